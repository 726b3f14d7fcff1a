"""MI355X (gfx950) backend of IonSolver's extended-LBM MHD time step.

The compute path is libionsolver_hip.so (hand-written HIP kernels behind the C
ABI of include/ionsolver.h); this package is the host-side mirror of the
reference's Lbm/LbmConfig API over that ABI, plus scenario setups.
"""
from .lbm import (Field, FieldMode, FieldSolver, FloatType, GraphicsConfig, IonError, Lbm, LbmConfig, ModelType,
                  Propellant, RelaxationTime, Units, VelocitySet, device_codec, hbm_copy_probe, hbm_pull_probe,
                  hbm_stream_probe, load_library, LIB_PATH, ABI)

__all__ = ["Field", "FieldMode", "FieldSolver", "FloatType", "GraphicsConfig", "IonError", "Lbm", "LbmConfig", "ModelType",
           "Propellant",
           "RelaxationTime", "Units", "VelocitySet", "device_codec", "hbm_copy_probe", "hbm_pull_probe",
           "hbm_stream_probe", "load_library", "LIB_PATH", "ABI"]
