"""MI355X-native BGK_PCM time step (thermal D2Q9 lattice Boltzmann with a
phase-change field), driven through the C ABI of include/bgk_pcm.h."""
from .params import Params, read_input_dat, write_input_dat, FP64, FP32_DEV, FP32_MOM, FP64_MOM, ROUND_STEP7, ROUND_CLEAN

__all__ = ["Params", "read_input_dat", "write_input_dat", "FP64", "FP32_DEV", "FP32_MOM", "FP64_MOM",
           "ROUND_STEP7", "ROUND_CLEAN"]
