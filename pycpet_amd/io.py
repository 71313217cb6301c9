"""Options, input parsing, atom filters and output writers.

Mirrors the hot-path parts of the reference's CPET/utils/io.py with the same
function names, arguments, return values and error behaviour:

  default_options_initializer / check_valid_options   io.py:200-266
  get_input_files                                      io.py:10-26
  filter_radius, filter_radius_whole_residue           io.py:269-338
  filter_atomspec, filter_in_box                       io.py:341-475
  parse_pqr, parse_pdb                                 io.py:478-636
  get_atoms_for_axes, calculate_center                 io.py:639-702
  save_numpy_as_dat                                    io.py:50-109

The radius and box filter decisions can also run on the GPU
(pycpet_amd.frame); the functions here keep the reference's numpy
semantics so both can be compared on identical coordinates.

PDB and PQR parsing and the .dat / .top writers run in the native library
(cpet_pdb_parse, cpet_pqr_parse, cpet_write_rows: host code, no GPU
needed), which reproduces the reference's parsed values and output bytes
exactly.
"""
from __future__ import annotations

import logging
from glob import glob

import numpy as np

log = logging.getLogger(__name__)

ATOM_ATTRIBUTES = ["atom_number", "atom_type", "resid", "resnum", "chain"]


def get_input_files(inputpath, filetype):
    if filetype not in ["pdb", "pqr"]:
        raise ValueError("filetype must be 'pdb' or 'pqr'")
    files_input = glob(inputpath + f"/*.{filetype}")
    if len(files_input) == 0:
        raise ValueError("No pdb files found in the input directory")
    if len(files_input) == 1:
        log.warning("Only one pdb file found in the input directory")
    return files_input


def default_options_initializer(options):
    """Fills missing option keys with the reference's defaults."""
    defaults = {
        "inputpath": "./inpdir",
        "outputpath": "./outdir",
        "inputfiletype": "pdb",
        "profile": False,
        "write_transformed_pdb": False,
        "dimensions": None,
        "step_size": None,
        "initializer": "uniform",
        "box_shift": [0, 0, 0],
        "filter_intersect": False,
        "concur_slip": 4,
        "dtype": "float32",
        "GPU_batch_freq": 100,
        "n_samples": None,
    }
    for k, v in defaults.items():
        if k not in options.keys():
            options[k] = v
    return options


def check_valid_options(options):
    if options["inputfiletype"] not in ["pdb", "pqr"]:
        raise ValueError("inputfiletype must be 'pdb' or 'pqr'")
    if options["initializer"] not in ["uniform", "random"]:
        raise ValueError("initializer must be 'uniform' or 'random'")
    if "x" in options.keys() and "y" in options.keys():
        if options["x"] == options["y"]:
            raise ValueError("x and y points used for vectors cannot be the same")


# ----------------------------------------------------------------- filters

def filter_radius(x, Q, center, radius=2.0):
    """Keeps charges with |x - center| < radius."""
    mask = np.linalg.norm(x - center, axis=1) < radius
    x_filtered = x[mask]
    Q_filtered = Q[mask]
    log.info("radius filter leaves: %d", len(Q_filtered))
    return x_filtered, Q_filtered


def residue_runs(ID):
    """Run index of each row: consecutive rows with equal (resid, resnum)
    share a run (a residue); equal ids that are not adjacent start new runs."""
    n = len(ID)
    if n == 0:
        return np.zeros(0, dtype=np.int32)
    resids = np.array([row[2] for row in ID])
    resnums = np.array([row[3] for row in ID])
    new = np.ones(n, dtype=bool)
    new[1:] = (resids[1:] != resids[:-1]) | (resnums[1:] != resnums[:-1])
    return (np.cumsum(new) - 1).astype(np.int32)


def filter_radius_whole_residue(x, Q, ID, center, radius=2.0, keep=None):
    """Keeps whole residues (runs of equal (resid, resnum)) that have at least
    one atom with |x - center| < radius.  `keep` may carry a precomputed
    keep-mask (e.g. from the GPU kernel) in place of the numpy decision."""
    if keep is None:
        inside = np.linalg.norm(x - center, axis=1) < radius
        runs = residue_runs(ID)
        any_in = np.zeros(runs.max() + 1 if len(runs) else 0, dtype=bool)
        np.logical_or.at(any_in, runs, inside)
        keep = any_in[runs]
    drop = np.flatnonzero(~keep)
    x_filtered = np.delete(x, drop, axis=0)
    Q_filtered = np.delete(Q, drop, axis=0)
    ID_filtered = np.delete(ID, drop, axis=0)
    log.info("radius filter leaves: %d", len(Q_filtered))
    return x_filtered, Q_filtered, ID_filtered


def filter_atomspec(x, Q, ID, filter_dict, intersect):
    """Removes atoms matching the filter: union of listed attribute values,
    or (intersect=True) an ordered intersection across keys with "*" and ""
    as wildcards."""
    considered = []
    for key in filter_dict.keys():
        if key not in ATOM_ATTRIBUTES:
            raise ValueError(f"Key {key} not recognized. Must be one of {ATOM_ATTRIBUTES}")
        considered.append(key)
    drop = set()
    if not intersect:
        for key in considered:
            vals = filter_dict[key]
            col = ATOM_ATTRIBUTES.index(key)
            for j in range(len(ID)):
                if ID[j][col] in vals:
                    drop.add(j)
    else:
        lens = [len(filter_dict[k]) for k in considered]
        if len(set(lens)) != 1:
            raise ValueError(
                "All filter lists must be the same length when using intersection filtering")
        for i in range(lens[0]):
            want = {k: filter_dict[k][i] for k in considered}
            for j in range(len(ID)):
                ok = True
                for k, v in want.items():
                    if v != "*" and v != "" and ID[j][ATOM_ATTRIBUTES.index(k)] != v:
                        ok = False
                        break
                if ok:
                    drop.add(j)
    drop = list(drop)
    x_filtered = np.delete(x, drop, axis=0)
    Q_filtered = np.delete(Q, drop, axis=0)
    ID_filtered = np.delete(ID, drop, axis=0)
    if len(x_filtered) != len(Q_filtered) or len(x_filtered) != len(ID_filtered):
        raise ValueError("Unexpected: filtered arrays have different lengths")
    log.info("Number of atoms filtered out: %d", len(drop))
    return x_filtered, Q_filtered, ID_filtered


def filter_in_box(x, Q, ID, center, dimensions, keep=None):
    """Removes charges strictly inside +-dimensions of (x - center).  Note the
    reference calls this on coordinates already centred and rotated, so
    `center` is subtracted a second time (kept as is)."""
    if keep is None:
        xr = x - center
        inside = np.ones(xr.shape[0], dtype=bool)
        for i in range(3):
            inside &= (xr[:, i] > -dimensions[i]) & (xr[:, i] < dimensions[i])
        keep = ~inside
    return x[keep], Q[keep], np.asarray(ID)[keep]


# ------------------------------------------------------------------ parsing

_PQR_WIDTH = 16  # bytes per atom / residue name slot of the native parser


def parse_pqr(path_to_pqr):
    """PQR parser (CPET/utils/io.py:478-571): returns (x (N,3), Q (N,1),
    atom_number, residue_name, residue_number, atom_type) with the
    reference's dtypes, its token logic for merged record names and its
    column shift for residue numbers above 999.  Parsing runs in
    cpet_pqr_parse; a record that parser does not read exactly as Python
    would (non-ASCII bytes, underscores in numbers, very long tokens) or that
    is malformed is re-read by the line-by-line restatement, which returns
    Python's result or raises Python's exception."""
    import ctypes

    from . import _native as nat

    lib = nat.load()
    path = str(path_to_pqr).encode()
    n = ctypes.c_int(0)
    if lib.cpet_pdb_scan(path, ctypes.byref(n)) != nat.CPET_OK:  # same record rule
        raise OSError(nat.last_error(lib))
    N = n.value
    if N == 0:
        return _parse_pqr_text(path_to_pqr)
    W = _PQR_WIDTH
    xyz = np.empty((N, 3), np.float64)
    q = np.empty(N, np.float64)
    anum = np.empty(N, np.int32)
    rnum = np.empty(N, np.int32)
    atype = np.zeros(W * N, np.uint8)
    rname = np.zeros(W * N, np.uint8)
    ptr = lambda a: a.ctypes.data
    if lib.cpet_pqr_parse(path, N, W, ptr(xyz), ptr(q), ptr(anum), ptr(rnum), ptr(atype),
                          ptr(rname)) != nat.CPET_OK:
        return _parse_pqr_text(path_to_pqr)
    return (xyz, q.reshape(-1, 1), anum.astype(np.int64), _fixed_str(rname, W),
            rnum.astype(np.int64), _fixed_str(atype, W))


def _parse_pqr_text(path_to_pqr):
    """Line-by-line restatement of the reference PQR parser (the fallback of
    parse_pqr, and its check)."""
    x, Q, atom_num, res_name, res_num, atom_type = [], [], [], [], [], []
    with open(path_to_pqr) as fh:
        lines = fh.readlines()
    for line in lines:
        if not (line.startswith("ATOM") or line.startswith("HETATM")):
            continue
        tok = line.split()
        shift = 0
        res_ind = 5
        merged = len(tok[0]) > 6  # "HETATM12345" without a separating space
        if merged:
            res_ind = 4
            atom_num.append(int(tok[0][6:]))
            res_name.append(tok[2])
            atom_type.append(tok[1])
        else:
            atom_num.append(int(tok[1]))
            res_name.append(tok[3])
            atom_type.append(tok[2])
        if len(tok[res_ind]) > 4:
            res_val = int(tok[res_ind - 1][1:])
        else:
            res_val = int(tok[res_ind])
        res_num.append(res_val)
        if res_val > 999:
            shift += int(np.log10(res_val) - 2)
        coords = [line[30 + shift:38 + shift], line[38 + shift:46 + shift],
                  line[46 + shift:54 + shift]]
        charge = line[54 + shift:61 + shift].strip()
        try:
            tempq = float(charge)
            temp = [float(c.strip()) for c in coords]
        except ValueError as exc:
            raise ValueError(
                "Charge or coordinates is not a useable number. Check pqr file "
                f"formatting for the following line: {line}") from exc
        x.append(temp)
        Q.append(tempq)
    return (np.array(x), np.array(Q).reshape(-1, 1), np.array(atom_num),
            np.array(res_name), np.array(res_num), np.array(atom_type))


def _parse_pdb_text(pdb_file_path, get_charges, dtype):
    """Line-by-line parser for files with non-ASCII records, where Python's
    character columns and the native parser's byte columns differ."""
    xyz, Q, atom_number, residue_name, residue_number, atom_type, chains = ([] for _ in range(7))
    with open(pdb_file_path, "r") as fh:
        for line in fh:
            if line.startswith("ATOM") or line.startswith("HETATM"):
                atom_number.append(int(line[6:11].strip()))
                atom_type.append(line[12:16].strip())
                residue_name.append(line[17:20].strip())
                chains.append(line[21])
                residue_number.append(int(line[22:26].strip()))
                xyz.append([float(line[30:38].strip()), float(line[38:46].strip()),
                            float(line[46:54].strip())])
                if get_charges:
                    Q.append(float(line[55:64].strip()))
    return (np.array(xyz, dtype=dtype).reshape(-1, 3), np.array(Q, dtype=dtype).reshape(-1, 1),
            np.array(atom_number, dtype=dtype), np.array(residue_name),
            np.array(residue_number), np.array(atom_type), np.array(chains))


def _fixed_str(buf: np.ndarray, width: int) -> np.ndarray:
    """NUL-padded ASCII byte records -> the unicode array np.array(list_of_str)
    gives (item width = the longest string): the bytes widened to UCS-4 code
    units and viewed as strings, which drop their trailing NULs as numpy's
    do."""
    rec = buf.reshape(-1, width)
    if rec.shape[0] == 0:
        return np.array([], dtype="U1")
    nz = rec != 0
    w = max(1, int((width - np.argmax(nz[:, ::-1], axis=1))[nz.any(axis=1)].max(initial=0)))
    return np.ascontiguousarray(rec[:, :w]).astype(np.uint32).view(f"U{w}").reshape(-1)


def parse_pdb(pdb_file_path, get_charges=False, float32=True):
    """Fixed-column PDB parser (CPET/utils/io.py:574-636): returns
    (xyz, [Q (N,1),] atom_number, residue_name, residue_number, atom_type,
    chains) with the reference's dtypes.  Parsing runs in cpet_pdb_parse;
    malformed numeric fields raise ValueError as int()/float() would."""
    import ctypes

    from . import _native as nat

    dtype = "float32" if float32 else "float64"
    lib = nat.load()
    path = str(pdb_file_path).encode()
    n = ctypes.c_int(0)
    rc = lib.cpet_pdb_scan(path, ctypes.byref(n))
    if rc != nat.CPET_OK:
        raise OSError(nat.last_error(lib))
    N = n.value
    if N == 0:  # np.array([]) of each list, as the reference builds them
        e = np.array([])
        return ((np.array([], dtype=dtype),) + ((np.array([], dtype=dtype).reshape(-1, 1),)
                if get_charges else ()) + (np.array([], dtype=dtype), e, e, e, e))
    xyz = np.empty((N, 3), np.float64)
    q = np.empty(N, np.float64)
    anum = np.empty(N, np.int32)
    rnum = np.empty(N, np.int32)
    atype = np.zeros(4 * N + 1, np.uint8)
    rname = np.zeros(3 * N + 1, np.uint8)
    chain = np.zeros(N + 1, np.uint8)
    ptr = lambda a: a.ctypes.data
    rc = lib.cpet_pdb_parse(path, N, int(bool(get_charges)), ptr(xyz), ptr(q), ptr(anum),
                            ptr(rnum), ptr(atype), ptr(rname), ptr(chain))
    if rc != nat.CPET_OK:
        msg = nat.last_error(lib)
        if msg.startswith("non-ASCII"):
            out = _parse_pdb_text(pdb_file_path, get_charges, dtype)
            return (out[0],) + ((out[1],) if get_charges else ()) + out[2:]
        raise ValueError(msg)
    rest = (anum.astype(dtype), _fixed_str(rname[:3 * N], 3), rnum.astype(np.int64),
            _fixed_str(atype[:4 * N], 4), _fixed_str(chain[:N], 1))
    base = (xyz.astype(dtype),)
    if get_charges:
        return base + (q.astype(dtype).reshape(-1, 1),) + rest
    return base + rest


def get_atoms_for_axes(x, atom_type, residue_number, chains, options, seltype="center"):
    sel = options[seltype]
    if "chains" in sel.keys():
        if type(sel["chains"]) != list:
            raise ValueError("chains must be a list")
        if len(sel["chains"]) != len(sel["atoms"]):
            raise ValueError("chains must be the same length as atoms")
        wanted = [(k, v, chain) for atom_dict, chain in zip(sel["atoms"], sel["chains"])
                  for k, v in atom_dict.items()]
        return [pos for atom_res in wanted for idx, pos in enumerate(x)
                if (atom_type[idx], residue_number[idx], chains[idx]) == atom_res]
    wanted = set((k, v) for atom_dict in sel["atoms"] for k, v in atom_dict.items())
    return [pos for (atype, rnum), pos in zip(zip(atom_type, residue_number), x)
            if (atype, rnum) in wanted]


def calculate_center(coordinates, method):
    if method == "mean":
        return np.mean(coordinates, axis=0)
    elif method == "first":
        return coordinates[0]
    elif method == "inverse":
        return 2 * np.mean(coordinates[1:], axis=0) - coordinates[0]


# ------------------------------------------------------------------ writers

_ROW_DTYPES = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.float16): 2}


def write_rows(name, data, fmt="%.3f", header=""):
    """numpy.savetxt(name, data, fmt) for fmt "%.3f" or "%.18e" (numpy's
    default), byte-identical, with the rows formatted by cpet_write_rows on
    all host cores; `header` is written verbatim first."""
    from . import _native as nat

    codes = {"%.3f": 0, "%.18e": 1}
    if fmt not in codes:
        raise ValueError(f"unsupported row format {fmt!r}")
    a = np.asarray(data)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    if a.ndim != 2:
        raise ValueError("Expected 1D or 2D array, got %dD array instead" % a.ndim)
    if a.dtype not in _ROW_DTYPES:
        a = a.astype(np.float64)
    a = np.ascontiguousarray(a)
    lib = nat.load()
    nat.check(lib.cpet_write_rows(str(name).encode(), header.encode() if header else None,
                                  a.ctypes.data, _ROW_DTYPES[a.dtype], a.shape[0],
                                  max(1, a.shape[1]), codes[fmt], 0), lib)


def save_numpy_as_dat(meta_data, volume, name):
    """Writes a field/ESP volume in the CPET .dat layout (CPET/utils/io.py:
    50-109): a 7-line header (sample density, box, frame, centre, basis) then
    rows formatted %.3f."""
    dims = meta_data["dimensions"]
    steps = meta_data["num_steps"]
    T = meta_data["transformation_matrix"].transpose()
    c = meta_data["center"]
    header = (
        "#Sample Density: {} {} {}; Volume: Box: {} {} {}\n".format(
            steps[0], steps[1], steps[2], dims[0], dims[1], dims[2])
        + "#Frame 0\n"
        + "#Center: {} {} {}\n".format(c[0], c[1], c[2])
        + "#Basis Matrix:\n# {} {} {}\n# {} {} {}\n# {} {} {}\n".format(
            T[0][0], T[0][1], T[0][2], T[1][0], T[1][1], T[1][2], T[2][0], T[2][1], T[2][2])
    )
    write_rows(name, volume, "%.3f", header)
