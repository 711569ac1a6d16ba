"""alien_amd: an MI355X-native (gfx950) Krylov backend with Alien's algebra API.

The product is libalien_hip.so (C++ host code + hand-written HIP kernels,
C ABI in include/alien_hip.h). This package builds it (`alien_amd.build`) and
binds its C ABI through ctypes (`alien_amd._capi`) for tests and benchmarks.
"""
from ._capi import (  # noqa: F401
    BICGS, CG, EXPORTED, GEN_CONVDIFF7, GEN_LOGNORMAL7, GEN_POISSON7, LAYOUT_BELLPACK, LAYOUT_CSR, LAYOUT_RBCSR,
    LAYOUT_RBDIA, MODE_FUSED, MODE_GRAPH, MODE_MIXED, MODE_REFERENCE, MODE_REFERENCE_FUTURE, PREC_CHEB, PREC_DIAG,
    PREC_ILU0,
    PREC_FILU0, PREC_LEARNED, PREC_NEUMANN, PREC_NONE, AlienHipError, Context, bellpack_bro, bellpack_build, check, gen_bsr2,
    gen_csr, gen_nnz,
    Matrix, gnn_forward, gnn_init, gnn_loss_grad, gnn_nparams, gnn_train, group_by_level, host_color, host_dcol,
    host_diagonals, host_levels, lib, mtx_read_csr, mtx_read_vector, mtx_write_csr, mtx_write_vector,
    rbcsr_bundle_offsets, rbdia_build, solve_csr, uniform01, vblock_to_csr)

__version__ = "0.1.0"
