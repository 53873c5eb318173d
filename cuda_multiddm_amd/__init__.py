"""MI355X-native multi-scale DDM engine (tile -> FFT -> lag difference ->
masked azimuthal average), drop-in for the analysis path of
zhangyuyi99/cuda_multiDDM.  The compute path is the HIP library built by
``cuda_multiddm_amd.build``; this package holds its Python front end."""
from .engine import (ACC_COMPENSATED, ACC_PLAIN, CLI_PATH, DDM_EBUSY, FFT_LDS, FFT_REGISTER, LIB_PATH, NORM_FRAMES, NORM_PAIRS, PIX_U8, PIX_U16BE, PIX_U16LE, PREC_BF16X2,
                     PREC_FP16X2, PREC_FP32, SPECTRUM_FULL, SPECTRUM_MASKED, TILER_FUSED, TILER_PER_SCALE, DDMError, Server,
                     Engine, k_fit, k_lag_accumulate, k_tile_fft, lib, lin_lambdas, log_lags, log_lambdas,
                     ring_mask, run_ddm, synth_brownian, synth_uniform, validate, write_iqt_file)

__all__ = [
    "CLI_PATH", "DDM_EBUSY", "Server", "LIB_PATH", "NORM_FRAMES", "NORM_PAIRS", "PIX_U8", "PIX_U16BE", "PIX_U16LE", "DDMError",
    "PREC_FP32", "PREC_BF16X2", "PREC_FP16X2", "ACC_COMPENSATED", "ACC_PLAIN", "SPECTRUM_FULL", "SPECTRUM_MASKED",
    "FFT_REGISTER", "FFT_LDS", "TILER_FUSED", "TILER_PER_SCALE",
    "Engine", "k_fit", "k_lag_accumulate", "k_tile_fft", "lib", "lin_lambdas", "log_lags", "log_lambdas", "ring_mask",
    "run_ddm", "synth_brownian", "synth_uniform", "validate", "write_iqt_file",
]
