"""MI355X-native Lambda_b -> J/psi p K- amplitude likelihood (HIP kernels behind a C ABI)."""
from . import abi, models  # noqa: F401
from ._lib import Context, PWAError, generate_phsp, generate_phsp_p4, helicity_couplings, lib, phsp_angles  # noqa: F401
