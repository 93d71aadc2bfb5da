"""MI355X-native Cholesky FGLS (drop-in for the reference's fgls_chol_gpu path).

Import name: ``cugwas_proto_amd`` (the directory name carries a hyphen, so
``load_package()`` in the repository's entry points registers it).
"""
from .fgls import (  # noqa: F401
    DeviceContext,
    EigenContext,
    FglsError,
    colmajor,
    device_count,
    exported_symbols,
    eigen_decompose,
    eigen_decompose_device,
    factor_M,
    fgls_chol_gpu,
    fgls_eigen,
    last_error,
    last_run_stats,
    lib,
    LIB_PATH,
    pack_i16,
    pack_u2,
    tile_i8,
    tile_u2,
    tile_u2_packed,
    to_i8,
    u2_to_i8,
    x2t_bytes,
    FGLS_OK,
    FGLS_ERR_ARG,
    FGLS_ERR_HIP,
    FGLS_ERR_NOMEM,
    FGLS_ERR_NOTSPD,
    FGLS_ERR_IO,
    FGLS_ERR_STATE,
    FGLS_MODE_FP64,
    FGLS_MODE_MIXED_I8,
    FGLS_SLICES_AUTO,
    FGLS_SLICES_SIX,
)
from .adjoint import SnpLoss, snp_loss  # noqa: F401
