"""MI355X-native FlashMatrix SpMM path (HIP for gfx950 behind a C ABI).

The public surface mirrors the reference's fm::sparse_matrix operator API;
see flashx_amd.fmx for the bindings.
"""
from .fmx import (Context, FmxError, SparseMatrix, compute_eigen, colnorm, default_context,  # noqa: F401
                  gemm_tall, gram, rmat_edges, row_scale)
