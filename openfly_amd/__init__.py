"""openfly_amd: MI355X-native saddle-search hot path of openFLY (neighbour lists, EAM/FS energy,
gradient and Hessian-vector products, L-BFGS and dimer) behind the reference's API, over the C-ABI of
libflygpu.so (include/flygpu.h)."""
from .flygpu import (  # noqa: F401
    DEVICE, EXIT_NAMES, FP64, HOST, MIXED, EAM, Box, Context, DataEAM, DeviceError, Dimer, FlyError, LBFGS, List,
    SplineOutOfBounds, StepLBFGS, Batch, Busy, Stopped, Server, batch_options, declared_symbols, dimer_options, lbfgs_options,
    lib, rotor_options, server_options, mech_options, MECH_NAMES, perturb, theta_tol, Rotor, log_prod_eigen, prefactor)
