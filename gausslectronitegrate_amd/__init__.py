"""MI355X (gfx950) integral engine for gausslet orbitals on cubic lattices.

The C ABI lives in lib/libgli_amd.so (declarations: include/*.h); this package
mirrors the reference Python module surface on top of it.
"""
from .pymodule import (  # noqa: F401
    grid_num_points, grid_point_to_linear_index, linear_index_to_grid_point,
    compute_kinetic_gausslet_integral, compute_kinetic_gausslet_integrals,
    compute_nuclear_gausslet_integral, compute_nuclear_gausslet_integrals,
    compute_eri_gausslet_integral, compute_sparse_eri_gausslet_integrals,
    compute_erida_gausslet_integral, compute_sparse_erida_gausslet_integrals,
    compute_eri_gausslet_integral_batch, compute_erida_gausslet_integral_batch,
    get_max_openmp_threads, SparseERI, SparseERIDA,
    load_sparse_eri_gausslet_integrals, save_sparse_eri_gausslet_integrals,
)
from ._lib import ERI_TENSOR_FP64, ERI_TENSOR_MIXED, ERI_DIRECT_FP64  # noqa: F401
