"""cpsz_amd: critical-point-preserving compression of 2D/3D float vector
fields on AMD Instinct MI355X (gfx950), stream-compatible with bzhan166/cpSZ.

The compute path is the HIP library ``libcpsz_hip.so`` (C ABI in
``include/cpsz.h``); this package is its Python binding.
"""
from .api import (CpszError, compress_2d_dev, compress_3d_dev, critical_points_2d_dev,  # noqa: F401
                  critical_points_3d_dev, decompress_2d_dev, decompress_3d_dev, device_count,
                  last_stages, lib, set_timing, stream_size_2d,
                  sz_compress_cp_preserve_2d_offline, sz_compress_cp_preserve_2d_offline_log,
                  sz_compress_cp_preserve_3d_offline_log, sz_decompress_cp_preserve_2d_offline,
                  sz_decompress_cp_preserve_2d_offline_log, sz_decompress_cp_preserve_3d_online_log,
                  sz_lossless_compress, sz_lossless_decompress)
from .api import batch_2d_config, batch_2d_set_config, compress_2d_batch_dev, decompress_2d_batch_dev  # noqa: F401
from .api import batch_3d_config, compress_3d_batch_dev, decompress_3d_batch_dev  # noqa: F401
from .api import compress_2d_learned_batch_dev  # noqa: F401
from .api import (compress_2d_learned_dev, gemm_bf16_dev, learn_analytic_eb_2d_dev,  # noqa: F401
                  learn_features_2d_dev, learn_init_params, learn_labels_2d_dev, learn_param_count,
                  learn_train_2d_dev, mlp_forward_dev, mlp_grad_dev, sz_compress_cp_preserve_2d_learned)
from .api import (compress_3d_learned_batch_dev, learn_analytic_eb_3d_dev, learn_features_3d_dev,  # noqa: F401
                  learn_init_params_3d, learn_labels_3d_dev, learn_param_count_3d, learn_train_3d_dev,
                  mlp_forward_3d_dev, mlp_grad_3d_dev, tet_cp_emulate)
from .api import check_fast_bounds  # noqa: F401
from .api import set_exact_2d_bounds  # noqa: F401
from .api import ratio_ge1_host, tri_bound_2d_fast_host  # noqa: F401
from .api import critical_points_2d_sos_dev, critical_points_3d_sos_dev, sos_det  # noqa: F401
from .api import sz_lossless_compress_batch_dev, sz_lossless_compress_dev  # noqa: F401
