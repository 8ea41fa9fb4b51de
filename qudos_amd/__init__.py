"""MI355X-native single-particle solve for QUDOS (8-band k.p plane-wave quantum dots).

The compute path is the C ABI of ``lib/libqdgpu.so`` (declared in ``include/qudos_gpu.h``);
this package is a thin ctypes binding over it plus a ``torch.autograd.Function`` for gradients.
"""
from ._lib import (QD_OK, QD_E_ARG, QD_E_INPUT, QD_E_ALLOC, QD_E_DEVICE, QD_E_NOCONV, QD_E_UNSUPPORTED, QD_E_NAN,
                   QD_FP64, QD_FP32_REFINE, QD_NTAB, QD_REDUCE_TWO_STAGE, QD_REDUCE_ONE_STAGE, QD_DP_FX, QD_DP_FY,
                   QD_DP_FZ, QD_DP_L1, QD_DP_L2, QD_DP_L3, QD_DP_L12, QD_DP_X, QdError, Inputs, Tables, Stats,
                   CoulombStats, AdjointStats, Solver, lib_path, load_library, read_inputs, build_tables, window,
                   write_coulomb, tables_derivative, tables_vjp, QD_DP_OX, QD_DP_OY, QD_DP_OZ, QD_DP_L4,
                   MomentumStats, selftest_adjoint_bin_host, selftest_momentum_host,
                   selftest_coulomb_upsample_host, selftest_hamiltonian_host, selftest_shape_host,
                   EXPORTED_SYMBOLS)

__all__ = ["QD_OK", "QD_E_ARG", "QD_E_INPUT", "QD_E_ALLOC", "QD_E_DEVICE", "QD_E_NOCONV", "QD_E_UNSUPPORTED",
           "QD_E_NAN", "QD_FP64", "QD_FP32_REFINE", "QD_NTAB", "QD_REDUCE_TWO_STAGE", "QD_REDUCE_ONE_STAGE", "QdError",
           "Inputs", "Tables", "Stats", "CoulombStats", "AdjointStats", "Solver", "lib_path", "load_library",
           "read_inputs", "build_tables", "window", "write_coulomb", "tables_derivative", "EXPORTED_SYMBOLS",
           "QD_DP_FX", "QD_DP_FY", "QD_DP_FZ", "QD_DP_L1", "QD_DP_L2", "QD_DP_L3", "QD_DP_L12", "QD_DP_X", "QD_DP_OX", "QD_DP_OY", "QD_DP_OZ", "QD_DP_L4", "tables_vjp",
           "MomentumStats", "selftest_adjoint_bin_host", "selftest_momentum_host",
           "selftest_coulomb_upsample_host", "selftest_hamiltonian_host", "selftest_shape_host"]
