"""vmps_amd: MI355X-native effective-Hamiltonian application (HxV) for
two-site DMRG pivots, behind the pivot API of TNWorkers/VMPS."""
from .pivot import PivotMatrix2, HxV, dot  # noqa: F401
