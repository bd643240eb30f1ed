"""skfemntv -- MI355X (gfx950) build of scikit-fem-native's assembly path.

Drop-in for ``skfemntv.asm(form, basis, **coefficients)`` with
``BilinearForm`` / ``LinearForm``, ``Basis`` over tetrahedral P1/P2 scalar
and vector fields, and the affine Nedelec-N1 ``TetN1Assembler``.  Assembly
runs in hand-written HIP kernels (libskfem_amd.so) reached through a C ABI;
PyTorch-ROCm only provides device buffers and the stream.
"""

from ._symbols import UnsupportedNativeForm
from .basis import Basis, GeometryDiagnostics
from .element import ElementComposite, ElementTetP1, ElementTetP2, ElementVector
from .forms import BilinearForm, Functional, LinearForm, asm
from ._assembly import (AssemblyMemoryBudgetError, AssemblyMemoryEstimate, LinearFormDiagnostics,
                        NativeBilinearForm, NativeCompositeBilinearForm, NativeCompositeLinearForm,
                        NativeCrossBilinearForm, NativeLinearForm)
from .hcurl import (AffineTetN1Basis, HcurlTetGeometryDiagnostics, TetN1Assembler,
                    estimate_tet_n1_assembly_memory)
from .mesh import MeshTet, MeshTet2
from .facet import FacetBasis, InteriorFacetBasis, NativeFacetBilinearForm, NativeFacetLinearForm
from .cut import CutBilinearFormAssembler, CutCrossAssembler, CutLinearFormAssembler
from .cut_basis import CutCellBasis, CutCellQuadrature, DiscreteField
from .runtime import (BACKEND, CAPABILITIES, available_num_threads, get_num_threads, has_capability,
                      set_num_threads, thread_limit)
from ._native import NativeUnavailable, NativeError

__version__ = "0.3.1+mi355x"

__all__ = [
    "__version__", "BACKEND", "CAPABILITIES", "available_num_threads", "get_num_threads", "has_capability",
    "set_num_threads", "thread_limit", "AssemblyMemoryBudgetError", "AssemblyMemoryEstimate",
    "AffineTetN1Basis", "HcurlTetGeometryDiagnostics", "TetN1Assembler", "estimate_tet_n1_assembly_memory",
    "LinearFormDiagnostics", "NativeLinearForm", "NativeBilinearForm", "NativeCrossBilinearForm",
    "NativeCompositeBilinearForm", "NativeCompositeLinearForm", "ElementComposite", "FacetBasis", "InteriorFacetBasis",
    "NativeFacetLinearForm", "NativeFacetBilinearForm",
    "CutLinearFormAssembler", "CutBilinearFormAssembler", "CutCrossAssembler", "CutCellBasis",
    "CutCellQuadrature", "DiscreteField",
    "LinearForm", "BilinearForm", "Functional", "Basis", "GeometryDiagnostics", "ElementTetP1", "ElementTetP2",
    "ElementVector", "MeshTet", "MeshTet2", "asm", "UnsupportedNativeForm", "NativeUnavailable", "NativeError",
]
