"""Options-file driver for the hot-path methods, mirroring the reference's
`CPET` class (CPET/source/CPET.py:24-229).

`CPET(options).run()` dispatches "topo", "topo_GPU", "volume",
"volume_ESP", "point_field" and "point_mag" with the reference's frame loop:
pick an unfinished input at random, skip it when its output already exists
(so several processes can share a directory and a rerun resumes), build a
`calculator`, compute, and write <frame>.top / <frame>_efield.dat /
<frame>_esp.dat / point_field.dat in the reference formats.  Methods outside
the hot path (cluster, pca, visualize, box_check) raise NotImplementedError.

"topo" / "topo_GPU" run up to `frames_per_batch` unfinished frames (an
option of this build, default 16; 1 gives the reference's one frame per
call) through one batched streamline launch and write their .top files as
soon as the launch's results are back.  The batch computes each frame
exactly as the per-frame call does, so the files are the same; frames that
alone cannot fill the GPU run together.  With the option integrator = "rk4"
(an extension; the reference integrates with Euler) frames run one per call.

Crash and sharing behaviour.  Each .top is written to a temporary name and
renamed into place, so a crash never leaves a partial file that a rerun
would take for a finished frame.  A crash loses the frames of the group in
flight: at most `frames_per_batch` frames (the reference, which writes after
every frame, CPET/source/CPET.py:119-123, loses at most one).  Several
processes may share one output directory as with the reference: before a
group is launched, frames whose .top appeared meanwhile are dropped from it,
so a frame is computed twice only when two processes launch it at the same
time, as in the reference.
"""
from __future__ import annotations

import logging
import os
from random import choice

import numpy as np

from .calculator import calculator
from .io import (check_valid_options, default_options_initializer, get_input_files,
                 save_numpy_as_dat, write_rows)

log = logging.getLogger(__name__)


def _write_atomic(path, write):
    """write(tmp) to a temporary file beside `path`, then rename it to
    `path` (atomic on one filesystem)."""
    tmp = f"{path}.part{os.getpid()}"
    try:
        write(tmp)
        os.replace(tmp, path)
    finally:
        if os.path.exists(tmp):
            os.unlink(tmp)


HOT_METHODS = ("topo", "topo_GPU", "volume", "volume_ESP", "point_field", "point_mag")


class CPET:
    def __init__(self, options):
        self.options = default_options_initializer(options)
        check_valid_options(self.options)
        self.log = log
        self.m = self.options["CPET_method"]
        self.inputpath = self.options["inputpath"]
        self.outputpath = self.options["outputpath"]
        self.inputfiletype = self.options["inputfiletype"]
        if not os.path.exists(self.outputpath):
            self.log.warning("Output directory does not exist, creating: %s", self.outputpath)
            os.makedirs(self.outputpath)

    def run(self):
        if self.m == "topo":
            self.run_topo()
        elif self.m == "topo_GPU":
            self.run_topo(gpu=True)
        elif self.m == "volume":
            self.run_volume()
        elif self.m == "volume_ESP":
            self.run_volume_ESP()
        elif self.m in ("point_field", "point_mag"):
            self.run_point()
        else:
            raise NotImplementedError(
                f"Method {self.m} is not one of the accelerated methods {HOT_METHODS}")

    def _frames(self, suffix):
        """Yields (file, frame) for unfinished inputs in random order."""
        files = get_input_files(self.inputpath, filetype=self.inputfiletype)
        while files:
            f = choice(files)
            files.remove(f)
            frame = f.split("/")[-1].split(".")[0]
            done = [p for p in os.listdir(self.outputpath) if p.endswith(suffix)]
            if frame + suffix in done:
                self.log.info("Already done for protein: %s, skipping...", frame)
                continue
            yield f, frame

    def run_topo(self, gpu=False):
        runtype = "compute_topo_GPU_batch_filter" if gpu else "compute_topo_complete_c_shared"
        per = max(1, int(self.options.get("frames_per_batch", 16)))
        group = []
        for f, frame in self._frames(".top"):
            group.append((frame, calculator(self.options, path_to_pdb=f)))
            if len(group) == per:
                self._topo_group(group, runtype)
                group = []
        if group:
            self._topo_group(group, runtype)

    def _topo_group(self, group, runtype):
        """Streamline features of the frames of `group` ((frame, calculator)
        pairs), one batched launch for several, then their .top files."""
        done = set(os.listdir(self.outputpath))
        for frame, _ in group:
            if frame + ".top" in done:
                self.log.info("Finished by another process meanwhile: %s, skipping...", frame)
        group = [(frame, c) for frame, c in group if frame + ".top" not in done]
        if not group:
            return
        euler = all(getattr(c, "integrator", "euler") == "euler" for _, c in group)
        batched = len(group) > 1 and euler  # the batch kernel integrates with Euler
        if batched:
            from .batch import batch_topo

            results = batch_topo([{"x": c.x, "q": c.Q, "starts": c.random_start_points,
                                   "n_iter": np.asarray(c.random_max_samples),
                                   "h": c.step_size, "dims": c.dimensions} for _, c in group])
            for (_, c), r in zip(group, results):
                c.topo_result = r
                c.hist = r.feats.astype(np.float32)
        for frame, c in group:
            self.calculator = c
            hist = c.hist if batched else getattr(c, runtype)()
            _write_atomic(self.outputpath + f"/{frame}.top",
                          lambda path: write_rows(path, hist, "%.18e"))

    def _meta(self, mesh_shape):
        c = self.calculator
        return {"dimensions": c.dimensions, "step_size": [c.step_size] * 3,
                "num_steps": [mesh_shape[0], mesh_shape[1], mesh_shape[2]],
                "transformation_matrix": c.transformation_matrix, "center": c.center}

    def run_volume(self):
        for f, frame in self._frames("_efield.dat"):
            self.calculator = calculator(self.options, path_to_pdb=f)
            field_box, mesh_shape = self.calculator.compute_box()
            meta = self._meta(mesh_shape)
            _write_atomic(self.outputpath + f"/{frame}_efield.dat",
                          lambda path: save_numpy_as_dat(name=path, volume=field_box,
                                                         meta_data=meta))

    def run_volume_ESP(self):
        for f, frame in self._frames("_esp.dat"):
            self.calculator = calculator(self.options, path_to_pdb=f)
            esp_box, mesh_shape = self.calculator.compute_box_ESP()
            meta = self._meta(mesh_shape)
            _write_atomic(self.outputpath + f"/{frame}_esp.dat",
                          lambda path: save_numpy_as_dat(name=path, volume=esp_box,
                                                         meta_data=meta))

    def run_point(self):
        files = get_input_files(self.inputpath, filetype=self.inputfiletype)
        if self.m == "point_field":
            outfile, runtype = self.outputpath + "/point_field.dat", "compute_point_field"
        else:
            outfile, runtype = self.outputpath + "/point_mag.dat", "compute_point_mag"
        with open(outfile, "w") as fh:
            for f in files:
                self.calculator = calculator(self.options, path_to_pdb=f)
                frame = f.split("/")[-1].split(".")[0]
                fh.write("{}:{}\n".format(frame, getattr(self.calculator, runtype)()))
