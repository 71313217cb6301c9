"""Writes the golden vectors in this directory from the reference's own
native kernels (CPET/utils/math_module.c compiled under oracle/_ref by
`make -C oracle ref`), driven the way the reference's ctypes wrappers drive
them (CPET/utils/c_ops.py:250-341 and CPET/utils/calculator.py:430-475).

Run in a checkout where the reference source exists:
    make -C oracle ref && python tests/golden/make_golden.py
    python tests/golden/make_golden.py pdb      # only the text fixtures
    python tests/golden/make_golden.py dipole   # only ref_topo_dipole.npz

Each .npz holds the seeded inputs and the reference's FP32 outputs only.
examples_point_field/ holds the reference's point-field example: its two
frames (examples/1A_point-field/pdb, gzipped), and in examples.json the
options files of examples 1A, 1B and 5A with the point_field.dat each wrote
(examples/*/outdir/point_field.dat).
ref_pdb_rows.pdb holds lines of the reference's own test frame
(tests/test_files/test_large.pdb): its header, first records, a TER line
and the last HETATM records (5-digit serials, 4-digit residue numbers).
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.oracle import Ref  # noqa: E402
from pycpet_amd import workloads as W  # noqa: E402

OUT = Path(__file__).resolve().parent


def volume(ref):
    x, q = W.ball_exclusion(500, 15.0, 1.0, 2.0, seed=11)
    axes = W.grid_axes((1.0, 1.0, 1.0), 0.25)  # 9^3
    g = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    E = ref.compute_looped_field(g, x, q)
    phi = ref.esp_points(g, x, q)
    pf = ref.calc_field(np.zeros(3, np.float32), x, q)
    np.savez_compressed(OUT / "ref_volume_500.npz", x=x, q=q, ax=axes[0], ay=axes[1],
                        az=axes[2], E=E, phi=phi, point_field=pf)


def volume_floor(ref):
    # a charge 5e-4 A from a grid point exercises the r^2 >= 1e-6 floor
    x, q = W.ball_exclusion(64, 6.0, 0.5, 1.0, seed=12)
    axes = W.grid_axes((0.5, 0.5, 0.5), 0.25)  # 5^3
    x = np.concatenate([x, np.array([[0.25, 0.0005, -0.25]], np.float32)])
    q = np.concatenate([q, np.array([0.7], np.float32)])
    g = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    E = ref.compute_looped_field(g, x, q)
    np.savez_compressed(OUT / "ref_volume_floor.npz", x=x, q=q, ax=axes[0], ay=axes[1],
                        az=axes[2], E=E)


def topo(ref):
    c = W.small_topo(N=200, L=160, cap=300, seed=13, dims=(0.6, 0.6, 0.6), h=0.01)
    ret = ref.topo(c.starts, c.n_iter, c.h, c.dims, c.x, c.q)
    np.savez_compressed(OUT / "ref_topo_200.npz", x=c.x, q=c.q, starts=c.starts,
                        n_iter=c.n_iter, h=np.float32(c.h), dims=c.dims, ret=ret)


def point_field_examples():
    import gzip
    import json
    import re

    ex = Path("/root/reference/examples")
    out = OUT / "examples_point_field"
    out.mkdir(exist_ok=True)
    for f in sorted((ex / "1A_point-field" / "pdb").glob("*.pdb")):
        # mtime=0 keeps the archive bytes reproducible
        with open(out / (f.name + ".gz"), "wb") as fh, \
                gzip.GzipFile(filename=f.name, mode="wb", fileobj=fh, mtime=0) as gz:
            gz.write(f.read_bytes())
    cases = {}
    for d in ("1A_point-field", "1B_point-field-oriented", "5A_filter-ids"):
        opts = json.loads((ex / d / "options" / "options.json").read_text())
        expected = {}
        for line in (ex / d / "outdir" / "point_field.dat").read_text().splitlines():
            name, vec = line.split(":", 1)
            expected[name] = [float(v) for v in re.findall(r"[-+0-9.eE]+", vec)]
        cases[d] = {"options": opts, "expected": expected}
    (out / "examples.json").write_text(json.dumps(cases, indent=1) + "\n")


def pdb_rows():
    src = Path("/root/reference/tests/test_files/test_large.pdb").read_text().splitlines(True)
    ter = next(i for i, ln in enumerate(src) if ln.startswith("TER"))
    keep = src[:14] + [src[ter]] + src[-6:]
    (OUT / "ref_pdb_rows.pdb").write_text("".join(keep))


def topo_dipole(ref):
    """thread_operation_dipole of the reference C on seeded dipoles."""
    rng = np.random.default_rng(21)
    N, L = 80, 160
    d = rng.standard_normal((N, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x = (d * rng.uniform(2.0, 7.0, (N, 1))).astype(np.float32)
    mu = rng.normal(0.0, 0.3, (N, 3)).astype(np.float32)
    dims = np.array([0.5, 0.45, 0.55], np.float32)
    starts = rng.uniform(-0.4, 0.4, (L, 3)).astype(np.float32)
    n_iter = rng.integers(1, 150, L).astype(np.int32)
    h = np.float32(0.01)
    feats = np.stack([ref.thread_operation_dipole(starts[i], n_iter[i], h, dims, x, mu)
                      for i in range(L)])
    np.savez_compressed(OUT / "ref_topo_dipole.npz", x=x, mu=mu, dims=dims, starts=starts,
                        n_iter=n_iter, h=h, feats=feats)


if __name__ == "__main__":
    if sys.argv[1:] == ["dipole"]:  # only the dipole vectors
        topo_dipole(Ref())
        sys.exit(0)
    pdb_rows()
    point_field_examples()
    if sys.argv[1:] == ["pdb"]:  # text fixtures only
        sys.exit(0)
    r = Ref()
    volume(r)
    volume_floor(r)
    topo(r)
    topo_dipole(r)
    for p in sorted(OUT.glob("*.npz")):
        print(p.name, p.stat().st_size)
