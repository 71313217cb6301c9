"""Writes synthetic frames in the fixed-column PDB layout the reference
parser reads (coordinates in columns 31-54, charge in 56-64)."""
import numpy as np


def write_pdb(path, x, q, resnames=None, resnums=None, names=None, chain="A"):
    n = len(q)
    resnames = resnames if resnames is not None else ["ALA"] * n
    resnums = resnums if resnums is not None else [1 + i // 5 for i in range(n)]
    names = names if names is not None else [f"C{i % 5}" for i in range(n)]
    with open(path, "w") as fh:
        fh.write("CRYST1   80.000   80.000   80.000  90.00  90.00  90.00               1\n")
        for i in range(n):
            fh.write(f"ATOM  {i + 1:5d} {names[i]:<4s} {resnames[i]:>3s} {chain}{resnums[i]:4d}    "
                     f"{x[i, 0]:8.3f}{x[i, 1]:8.3f}{x[i, 2]:8.3f} {q[i]:8.4f}  1.8240       C  \n")
        fh.write("END\n")


def synthetic_frame(path, n=600, seed=0, center=(10.0, -4.0, 7.5), R=14.0):
    """A frame of n charges in residues of 5 atoms around `center`, with a
    marker residue (LIG 900: atoms C7, O9, N8) at the centre for the axes."""
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x = np.asarray(center) + d * (3.0 + (R - 3.0) * rng.random(n)[:, None])
    q = rng.normal(0, 0.4, n)
    q -= q.mean()
    x = np.round(x, 3)
    q = np.round(q, 4)
    names = [f"C{i % 5}" for i in range(n)]
    resnames = ["ALA" if (i // 5) % 2 else "GLY" for i in range(n)]
    resnums = [1 + i // 5 for i in range(n)]
    lig = np.asarray(center) + np.array([[0.6, 0.0, 0.0], [-0.6, 0.1, 0.0], [0.0, 0.9, 0.3]])
    x = np.concatenate([x, np.round(lig, 3)])
    q = np.concatenate([q, [0.1, -0.2, 0.1]])
    names += ["C7", "O9", "N8"]
    resnames += ["LIG"] * 3
    resnums += [900] * 3
    write_pdb(path, x, q, resnames, resnums, names)
    return x, q
