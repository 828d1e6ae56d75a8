"""MI355X (gfx950) batched reduced-Wong-Wang simulation: Python front end of
libbnm_hip.so.  See include/bnm_hip.h for the C ABI."""
from ._lib import (BNM_E_ARG, BNM_E_HIP, BNM_E_NODEV, BNM_E_SIZE, BNM_E_UNSUP, BNM_OK, CMAES, EXPORTS,
                   FC_BF16, FC_BF16X3, FC_F32, FC_F64, KERNEL_AUTO, KERNEL_FAST, KERNEL_MFMA, KERNEL_PLAIN,
                   LIB_PATH, MLP, Multi, PREC_BF16, PREC_BF16X3, PREC_F32, PREC_F64, BnmError, Config, Context,
                   analytical_fic, default_config, finalize, lib, make_noise, make_noise_chunked, n_pairs, regional_params, output_ts, run_simulations, sparse_layout,
                   window_bounds)
from . import workloads

__all__ = [
    "BNM_E_ARG", "BNM_E_HIP", "BNM_E_NODEV", "BNM_E_SIZE", "BNM_E_UNSUP", "BNM_OK", "CMAES", "EXPORTS",
    "FC_BF16", "FC_BF16X3", "FC_F32", "FC_F64", "KERNEL_AUTO", "KERNEL_FAST", "KERNEL_MFMA", "KERNEL_PLAIN",
    "LIB_PATH", "MLP", "Multi", "PREC_BF16", "PREC_BF16X3", "PREC_F32", "PREC_F64", "BnmError", "Config", "Context",
    "analytical_fic", "default_config", "finalize", "lib", "make_noise", "make_noise_chunked", "regional_params", "n_pairs", "output_ts", "run_simulations", "sparse_layout",
    "window_bounds", "workloads",
]
