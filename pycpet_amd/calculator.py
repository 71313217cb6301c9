"""Drop-in for the reference's per-frame `calculator`
(CPET/source/calculator.py:36-978).

Same constructor `calculator(options, path_to_pdb)`, same attributes (x, Q,
ID, center, mesh, transformation_matrix, random_start_points,
random_max_samples, step_size, dimensions, max_steps, ...), same compute
methods and return types.  The compute methods read those attributes at call
time, so scripts that overwrite `calc.x` / `calc.Q` between calls
(CPET/source/scripts/residue_breakdown_analysis.py:177-180) keep working.

Differences from the reference, all on the compute side:
  * compute_box / compute_box_ESP run one GPU launch over the whole grid
    (the reference loops in C, and for the ESP calls the C kernel once per
    grid point from Python);
  * compute_topo_complete_c_shared and compute_topo_GPU_batch_filter both run
    the streamline kernels over all starts in one launch and return rows in
    input order (the reference's GPU path returns them in dump order);
  * the radius / in-box filter decisions and (optionally) the box-frame
    transform run on the GPU (option "frame_backend": "gpu", the default, or
    "numpy").
"""
from __future__ import annotations

import logging
import time
from copy import deepcopy

import numpy as np

from . import field_ops
from . import frame as gpu_frame
from . import ops
from .io import (calculate_center, default_options_initializer, filter_atomspec,
                 filter_in_box, filter_radius_whole_residue, get_atoms_for_axes,
                 parse_pdb, parse_pqr, residue_runs)

log = logging.getLogger(__name__)


class calculator:
    def __init__(self, options, path_to_pdb=None):
        options = default_options_initializer(options)
        self.profile = options["profile"]
        self.log = log
        self.path_to_pdb = path_to_pdb
        self.frame_backend = options.get("frame_backend", "gpu")
        self.precision = options.get("precision", "auto")
        self.integrator = options.get("integrator", "euler")  # "rk4": extension

        self.step_size = options["step_size"]
        self.dimensions = np.array(options["dimensions"]) if "dimensions" in options.keys() else None
        self.n_samples = options["n_samples"] if "n_samples" in options.keys() else None
        self.concur_slip = options["concur_slip"]
        self.GPU_batch_freq = options["GPU_batch_freq"]
        self.dtype = options["dtype"]
        self.max_streamline_init = options.get("max_streamline_init", "true_rand")
        self.write_transformed_pdb = options["write_transformed_pdb"]
        self.strip_filter = options.get("strip_filter", False)
        self.box_shift = options.get("box_shift", [0, 0, 0])

        if ".pqr" in self.path_to_pdb:
            (self.x, self.Q, self.atom_number, self.resids, self.residue_number,
             self.atom_type) = parse_pqr(self.path_to_pdb)
        else:
            (self.x, self.Q, self.atom_number, self.resids, self.residue_number,
             self.atom_type, self.chains) = parse_pdb(self.path_to_pdb, get_charges=True)

        # ID rows (atom number, name, residue name, residue number[, chain]) of
        # the parsed arrays; built on first use (the filters), see ID below
        if hasattr(self, "chains"):
            self._id_cols = (self.atom_number, self.atom_type, self.resids,
                             self.residue_number, self.chains)
        else:
            self._id_cols = (self.atom_number, self.atom_type, self.resids,
                             self.residue_number)
            self.chains = [None] * len(self.x)
        self._ID = None

        # centre and axes (CPET/source/calculator.py:152-206)
        if type(options["center"]) == list:
            self.center = np.array(options["center"])
        elif type(options["center"]) == dict:
            pos = get_atoms_for_axes(self.x, self.atom_type, self.residue_number, self.chains,
                                     options, seltype="center")
            self.center = calculate_center(pos, method=options["center"]["method"])
        else:
            raise ValueError("option 'center' takes a coordinate list or an atom selection "
                             f"dict, not {type(options['center']).__name__}")
        if "x" not in options.keys():
            self.log.info("option 'x' absent: the box keeps the input frame's axes and 'y' is not read")
            compute_y = False
        elif type(options["x"]) == list:
            self.x_vec_pt = np.array(options["x"])
            compute_y = True
        elif type(options["x"]) == dict:
            pos = get_atoms_for_axes(self.x, self.atom_type, self.residue_number, self.chains,
                                     options, seltype="x")
            self.x_vec_pt = calculate_center(pos, method=options["x"]["method"])
            compute_y = True
        else:
            raise ValueError("option 'x' takes a coordinate list or an atom selection dict, "
                             f"not {type(options['x']).__name__}")
        if compute_y:
            if type(options["y"]) == list:
                self.y_vec_pt = np.array(options["y"])
            elif type(options["y"]) == dict:
                pos = get_atoms_for_axes(self.x, self.atom_type, self.residue_number,
                                         self.chains, options, seltype="y")
                self.y_vec_pt = calculate_center(pos, method=options["y"]["method"])
            else:
                raise ValueError("option 'y' must be a coordinate list or an atom selection dict "
                                 "whenever 'x' is given")

        self.x_copy = deepcopy(self.x)
        self.residue_number_copy = deepcopy(self.residue_number)
        self.resids_copy = deepcopy(self.resids)
        self.atom_number_copy = deepcopy(self.atom_number)
        self.atom_type_copy = deepcopy(self.atom_type)
        self.chains_copy = deepcopy(self.chains)

        if "filter" in options.keys():
            self.x, self.Q, self.ID = filter_atomspec(self.x, self.Q, self.ID, options["filter"],
                                                     options["filter_intersect"])
            if all(c is None for c in self.chains):
                del self.chains
            self.atom_number = [self.ID[i][0] for i in range(len(self.x))]
            self.atom_type = [self.ID[i][1] for i in range(len(self.x))]
            self.resids = [self.ID[i][2] for i in range(len(self.x))]
            self.residue_number = [self.ID[i][3] for i in range(len(self.x))]
            if hasattr(self, "chains"):
                self.chains = [self.ID[i][4] for i in range(len(self.x))]

        if "filter_radius" in options.keys():
            radius = float(options["filter_radius"])
            keep = None
            if self.frame_backend == "gpu" and len(self.x):
                keep = gpu_frame.radius_residue_keep(self.x, self.center, radius,
                                                     residue_runs(self.ID))
            self.x, self.Q, self.ID = filter_radius_whole_residue(
                x=self.x, Q=self.Q, ID=self.ID, center=self.center, radius=radius, keep=keep)

        assert "CPET_method" in options.keys(), "CPET_method must be specified"
        method = options["CPET_method"]
        if method in ("volume", "volume_ESP") and hasattr(self, "y_vec_pt"):
            N_cr = 2 * self.dimensions / self.step_size
            N_cr = [int(N_cr[0]), int(N_cr[1]), int(N_cr[2])]
            self.mesh, self.transformation_matrix = field_ops.initialize_box_points_uniform(
                center=self.center, x=self.x_vec_pt, y=self.y_vec_pt, N_cr=N_cr,
                dimensions=self.dimensions, dtype=self.dtype, inclusive=True)
        elif method in ("topo", "topo_GPU") and hasattr(self, "y_vec_pt"):
            self.max_steps = round(2 * np.linalg.norm(self.dimensions) / self.step_size)
            if options["initializer"] == "random":
                (self.random_start_points, self.random_max_samples, self.transformation_matrix,
                 self.max_streamline_len) = field_ops.initialize_box_points_random(
                    self.center, self.x_vec_pt, self.y_vec_pt, self.dimensions, self.n_samples,
                    dtype=self.dtype, max_steps=self.max_steps)
            elif options["initializer"] == "uniform":
                num_per_dim = round(self.n_samples ** (1 / 3))
                if num_per_dim ** 3 < self.n_samples:
                    num_per_dim += 1
                self.n_samples = num_per_dim ** 3
                seed = 42 if self.max_streamline_init == "fixed_rand" else None
                (self.random_start_points, self.random_max_samples,
                 self.transformation_matrix) = field_ops.initialize_box_points_uniform(
                    center=self.center, x=self.x_vec_pt, y=self.y_vec_pt,
                    dimensions=self.dimensions, N_cr=[num_per_dim] * 3, dtype=self.dtype,
                    max_steps=self.max_steps, ret_rand_max=True, inclusive=False, seed=seed)
                self.random_start_points = self.random_start_points.reshape(-1, 3)
                self.n_samples = len(self.random_start_points)
        elif method in ("point_field", "point_mag", "box_check") and hasattr(self, "y_vec_pt"):
            _, _, self.transformation_matrix = field_ops.initialize_box_points_uniform(
                center=self.center, x=self.x_vec_pt, y=self.y_vec_pt, dimensions=[0, 0, 0],
                N_cr=[0, 0, 0], dtype=self.dtype, max_steps=0, ret_rand_max=True,
                inclusive=False)

        # box frame (CPET/source/calculator.py:341-371)
        shift = None if self.box_shift == [0, 0, 0] else np.array(self.box_shift)
        Tinv = np.linalg.inv(self.transformation_matrix) if hasattr(self, "y_vec_pt") else None
        if self.frame_backend == "gpu" and len(self.x):
            # in-box filtering happens between rotation and shift
            self.x = gpu_frame.frame_transform(self.x, self.center, Tinv, None)
        else:
            self.x = (self.x - self.center) @ Tinv if Tinv is not None else self.x - self.center
        if "filter_in_box" in options.keys() and bool(options["filter_in_box"]):
            keep = None
            if self.frame_backend == "gpu" and len(self.x):
                keep = gpu_frame.in_box_keep(self.x, self.center, self.dimensions)
            self.x, self.Q, self.ID = filter_in_box(x=self.x, Q=self.Q, ID=self.ID,
                                                    center=self.center,
                                                    dimensions=self.dimensions, keep=keep)
        if shift is not None:
            self.x = self.x - shift

        if self.write_transformed_pdb:
            self._write_transformed_pdb(path_to_pdb, Tinv)

        if self.dtype == "float32":
            self.x = self.x.astype(np.float32)
            self.Q = self.Q.astype(np.float32)
            self.center = self.center.astype(np.float32)
            if self.dimensions is not None:
                self.dimensions = self.dimensions.astype(np.float32)
        self.log.info("Initialized calculator: %d charges", len(self.Q))

    @property
    def ID(self):
        """List of per-atom ID tuples, as the reference's constructor builds it
        from the parsed columns (CPET/source/calculator.py:125-145); a
        filter replaces it with its own result.  Built when first read (zip
        yields the same numpy scalars as indexing the columns), so frames
        that no filter touches skip it."""
        if getattr(self, "_ID", None) is None:
            cols = getattr(self, "_id_cols", None)
            if cols is None:  # made with __new__ and never given an ID
                raise AttributeError("'calculator' object has no attribute 'ID'")
            self._ID = list(zip(*cols))
        return self._ID

    @ID.setter
    def ID(self, value):
        self._ID = value

    def _write_transformed_pdb(self, path_to_pdb, Tinv):
        if self.strip_filter:
            xc = self.x
            nums, names, resids, resnums = (self.atom_number, self.atom_type, self.resids,
                                            self.residue_number)
        else:
            xc = (self.x_copy - self.center) @ Tinv if Tinv is not None else self.x_copy - self.center
            xc = xc + np.array(self.box_shift)
            nums, names, resids, resnums = (self.atom_number_copy, self.atom_type_copy,
                                            self.resids_copy, self.residue_number_copy)
        name = f"transform_{path_to_pdb.split('/')[-1][:-4]}.pdb"
        with open(name, "w") as fh:
            for i in range(len(xc)):
                x, y, z = xc[i]
                fh.write(f"ATOM  {int(float(nums[i])):>5d} {names[i]:<4} {resids[i]:>3} A"
                         f"{int(float(resnums[i])):>4d}    {x:8.3f}{y:8.3f}{z:8.3f}  1.00  0.00"
                         "           \n")

    # ------------------------------------------------------------- compute

    def compute_point_field(self):
        """Field at the box centre (the origin of the transformed frame)."""
        t0 = time.time()
        E = field_ops.calculate_electric_field_c_shared_full_alt(
            np.array([0, 0, 0]), self.x, self.Q)
        self.log.info("point field: %.6f s", time.time() - t0)
        return E

    def compute_point_mag(self):
        """|E| at the box centre.  (The reference calls a nonexistent np.norm
        here, CPET/source/calculator.py:433; this returns the magnitude.)"""
        return np.linalg.norm(self.compute_point_field())

    def compute_box(self):
        field_box = field_ops.compute_field_on_grid(self.mesh, self.x, self.Q,
                                                    prec=self.precision)
        return field_box, self.mesh.shape

    def compute_box_ESP(self):
        esp_box = field_ops.compute_ESP_on_grid(self.mesh, self.x, self.Q, prec=self.precision)
        return esp_box, self.mesh.shape

    def _topo(self):
        t0 = time.time()
        r = ops.topo(self.random_start_points, np.asarray(self.random_max_samples),
                     self.step_size, self.dimensions, self.x, self.Q,
                     integrator=getattr(self, "integrator", "euler"))
        self.topo_result = r
        self.hist = r.feats.astype(np.float32)
        self.log.debug("Time taken for %d streamlines with N_charges = %d: %.3f s",
                       len(self.random_start_points), len(self.Q), time.time() - t0)
        return self.hist

    def compute_topo_complete_c_shared(self):
        """(L,2) float32 [distance, mean curvature] per start, input order."""
        return self._topo()

    def compute_topo_GPU_batch_filter(self):
        """Same kernels as compute_topo_complete_c_shared (the reference's
        torch path computes the same quantities, CPET/utils/gpu.py)."""
        return self._topo()

    def compute_topo_complete_c_shared_dipole(self):
        """(L,2) float32 [distance, mean curvature] per start in the field of
        point dipoles self.x (N,3) with moments self.mu (N,3), read at call
        time like the reference (CPET/source/calculator.py:714-749, called by
        scripts/field_topology_dipole.py); the reference's dipole pair term
        (math_module.c:335-371, 593-659)."""
        t0 = time.time()
        r = ops.topo_dipole(self.random_start_points, np.asarray(self.random_max_samples),
                            self.step_size, self.dimensions, self.x, self.mu)
        self.topo_result = r
        self.hist = r.feats.astype(np.float32)
        self.log.debug("Time taken for %d calculations with N_dipoles = %d: %.3f s",
                       len(self.random_start_points), len(self.mu), time.time() - t0)
        return self.hist
