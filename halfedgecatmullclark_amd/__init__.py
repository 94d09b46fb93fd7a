"""MI355X-native uniform Catmull-Clark refinement of halfedge cages.

The product is the C-ABI library ``libcchip.so`` (declared in
``include/CatmullClarkHIP.h``), a drop-in for the ``ccs_Refine_*`` path of
CatmullClark.h backed by hand-written HIP kernels for gfx950.  This package
holds its sources (``csrc/``), the host-side cage tooling (``cage``), the
ctypes front end (``capi``) and a torch autograd wrapper (``autograd``).
"""
from . import cage  # noqa: F401
