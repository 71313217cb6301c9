"""Hot-path helpers with the names and semantics of the reference's
CPET/utils/calculator.py:

  initialize_box_points_random      :123-168  (same np.random call sequence)
  initialize_box_points_uniform     :171-248
  compute_field_on_grid             :430-447  -> one GPU launch (M,6) float32
  compute_ESP_on_grid               :450-475  -> one GPU launch (M,4) float16
  calculate_electric_field_c_shared_full_alt  :296-310  point field
  calculate_esp_c_shared_full       :330-341
  calculate_thread_c_shared         :344-348
  curv, compute_curv_and_dist       :501-561  (numpy, for inspection)
  Inside_Box, inside_box_mask       :564-593
"""
from __future__ import annotations

import logging

import numpy as np

from . import ops

log = logging.getLogger(__name__)


def _basis(center, x, y):
    x = x - center
    y = y - center
    x_unit = x / np.linalg.norm(x)
    y_unit = y / np.linalg.norm(y)
    z_unit = np.cross(x_unit, y_unit)
    z_unit = z_unit / np.linalg.norm(z_unit)
    y_unit = np.cross(z_unit, x_unit)
    y_unit = y_unit / np.linalg.norm(y_unit)
    return x_unit, y_unit, z_unit


def initialize_box_points_random(center, x, y, dimensions, n_samples, max_steps=10,
                                 dtype="float32"):
    """Uniform random starts in the box (local frame) and random step limits,
    drawn with the same global np.random calls in the same order as the
    reference, so a seeded global state reproduces its starts."""
    half_length, half_width, half_height = dimensions
    x_unit, y_unit, z_unit = _basis(center, x, y)
    random_x = np.random.uniform(-half_length, half_length, n_samples)
    random_y = np.random.uniform(-half_width, half_width, n_samples)
    random_z = np.random.uniform(-half_height, half_height, n_samples)
    pts = np.column_stack([random_x, random_y, random_z])
    if dtype == "float32":
        pts = pts.astype(np.float32)
    transformation_matrix = np.column_stack([x_unit, y_unit, z_unit]).T
    random_max_samples = np.random.randint(1, max_steps, n_samples)
    return pts, random_max_samples, transformation_matrix, max_steps


def initialize_box_points_uniform(center, x, y, N_cr, dimensions, dtype="float32",
                                  max_steps=10, ret_rand_max=False, inclusive=True, seed=None):
    """Tensor-product grid over the box: inclusive linspace(-d, d, N+1) per
    axis, or the N interior points of linspace(-d, d, N+1, endpoint=False)."""
    half_length, half_width, half_height = dimensions
    x_unit, y_unit, z_unit = _basis(center, x, y)
    transformation_matrix = np.column_stack([x_unit, y_unit, z_unit]).T
    if inclusive:
        xs = np.linspace(-half_length, half_length, N_cr[0] + 1)
        ys = np.linspace(-half_width, half_width, N_cr[1] + 1)
        zs = np.linspace(-half_height, half_height, N_cr[2] + 1)
    else:
        xs = np.linspace(-half_length, half_length, N_cr[0] + 1, endpoint=False)[1:]
        ys = np.linspace(-half_width, half_width, N_cr[1] + 1, endpoint=False)[1:]
        zs = np.linspace(-half_height, half_height, N_cr[2] + 1, endpoint=False)[1:]
    mx, my, mz = np.meshgrid(xs, ys, zs, indexing="ij")
    local_coords = np.stack([mx, my, mz], axis=-1, dtype=dtype)
    if seed is not None:
        np.random.seed(seed)
    if ret_rand_max:
        n = local_coords.shape[0] * local_coords.shape[1] * local_coords.shape[2]
        random_max_samples = np.random.randint(1, max_steps, n)
        return local_coords, random_max_samples, transformation_matrix
    return local_coords, transformation_matrix


def mesh_axes(grid_coords):
    """(ax, ay, az) if grid_coords (nx,ny,nz,3) is a tensor-product mesh in
    'ij' order, else None."""
    g = np.asarray(grid_coords)
    if g.ndim != 4 or g.shape[-1] != 3:
        return None
    ax, ay, az = g[:, 0, 0, 0], g[0, :, 0, 1], g[0, 0, :, 2]
    if not ((g[..., 0] == ax[:, None, None]).all() and (g[..., 1] == ay[None, :, None]).all()
            and (g[..., 2] == az[None, None, :]).all()):
        return None
    return ax.astype(np.float32), ay.astype(np.float32), az.astype(np.float32)


def compute_field_on_grid(grid_coords, x, Q, prec: str = "auto"):
    """(M,6) float32: grid point and softened field, one GPU launch."""
    x_0 = np.asarray(grid_coords).reshape(-1, 3)
    axes = mesh_axes(grid_coords)
    if axes is not None:
        E, _ = ops.field_grid(*axes, x, Q, want_E=True, prec=prec)
    else:
        E, _ = ops.field_points(x_0, x, Q, want_E=True, prec=prec)
    return np.concatenate((x_0, E), axis=1)


def compute_ESP_on_grid(grid_coords, x, Q, prec: str = "auto"):
    """(M,4) float16: grid point and potential (cast to half precision as the
    reference does before writing, CPET/utils/calculator.py:475)."""
    x_0 = np.asarray(grid_coords).reshape(-1, 3)
    axes = mesh_axes(grid_coords)
    if axes is not None:
        _, phi = ops.field_grid(*axes, x, Q, want_E=False, want_phi=True, prec=prec)
    else:
        _, phi = ops.field_points(x_0, x, Q, want_E=False, want_phi=True, prec=prec)
    return np.concatenate((x_0, phi.reshape(-1, 1)), axis=1).astype(np.half)


def calculate_electric_field_c_shared_full_alt(x_0, x, Q):
    """Unsoftened field at one point (calc_field semantics), float32 (3,)."""
    E, _ = ops.field_points(np.asarray(x_0, dtype=np.float32).reshape(1, 3), x, Q,
                            want_E=True, soften=False)
    return E[0]


def calculate_esp_c_shared_full(x_0, x, Q):
    _, phi = ops.field_points(np.asarray(x_0, dtype=np.float32).reshape(1, 3), x, Q,
                              want_E=False, want_phi=True)
    return phi.astype(np.float32)


def calculate_thread_c_shared(x_0, n_iter, x, Q, step_size, dimensions):
    r = ops.topo(np.asarray(x_0, dtype=np.float32).reshape(1, 3), np.array([n_iter]),
                 step_size, dimensions, x, Q)
    return r.feats[0].astype(np.float32)


def curv(v_prime, v_prime_prime, eps=10e-6):
    denominator = np.linalg.norm(v_prime, axis=0) ** 3
    curvature = np.linalg.norm(np.cross(v_prime, v_prime_prime), axis=0)
    return curvature / eps if denominator == 0 else curvature / denominator


def compute_curv_and_dist(x_init, x_init_plus, x_init_plus_plus, x_0, x_0_plus, x_0_plus_plus):
    curv_init = curv(x_init_plus - x_init, x_init_plus_plus - 2 * x_init_plus + x_init)
    curv_final = curv(x_0_plus - x_0, x_0_plus_plus - 2 * x_0_plus + x_0)
    return [np.linalg.norm(x_init - x_0, axis=-1), (curv_init + curv_final) / 2]


def inside_box_mask(points, dimensions):
    hl, hw, hh = dimensions
    return ((points[:, 0] >= -hl) & (points[:, 0] <= hl) & (points[:, 1] >= -hw)
            & (points[:, 1] <= hw) & (points[:, 2] >= -hh) & (points[:, 2] <= hh))


def Inside_Box(local_point, dimensions):
    hl, hw, hh = dimensions
    return (-hl <= local_point[0] <= hl and -hw <= local_point[1] <= hw
            and -hh <= local_point[2] <= hh)


def histogram_edges(lo: float, hi: float, nbins: int) -> np.ndarray:
    """Bin edges as numpy.histogram2d builds them for an explicit range."""
    if lo == hi:
        lo, hi = lo - 0.5, hi + 0.5
    return np.linspace(lo, hi, nbins + 1)
