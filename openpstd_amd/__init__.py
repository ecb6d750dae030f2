"""openpstd_amd — MI355X-native PSTD update loop behind openPSTD's KernelInterface.

The configuration types mirror kernel/KernelInterface.h; PSTDKernel(GPU=True)
runs the HIP solver in libopenpstd_amd.so (built by ``python -m openpstd_amd.build``).
"""
from .config import (PSTDSettings, DomainConfEdge, DomainConf, PSTDConfiguration, RK_COEFFICIENTS,
                     PSTD_DOMAIN_SIDE_TOP, PSTD_DOMAIN_SIDE_BOTTOM, PSTD_DOMAIN_SIDE_LEFT,
                     PSTD_DOMAIN_SIDE_RIGHT, PSTD_DOMAIN_SIDE_ALL, PSTD_DOMAIN_SIDE_NONE)
from .kernel import (PSTDKernel, KernelInterface, KernelCallback, KernelCallbackLog, SimulationMetadata,
                     PSTDKernelNotConfiguredException, DeviceSolver, BatchSolver, RecordingCallback)
from ._lib import PSTDError
from .pstd_file import PSTDFile

__all__ = ["PSTDSettings", "DomainConfEdge", "DomainConf", "PSTDConfiguration", "RK_COEFFICIENTS", "PSTDKernel",
           "KernelInterface", "KernelCallback", "KernelCallbackLog", "SimulationMetadata",
           "PSTDKernelNotConfiguredException", "DeviceSolver", "BatchSolver", "RecordingCallback", "PSTDError",
           "PSTDFile"]
