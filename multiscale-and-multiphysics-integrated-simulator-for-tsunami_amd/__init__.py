"""MI355X-native CADMAS-SURF/3D single-phase VOF-SMAC time step.

Host side of the drop-in: the reference's input deck (deck.py), its staggered
porous field layout (layout.py), its .rsl/.res and list-file per-step output
(rsl.py), its time loop (timeloop.py, driver.py), and the ctypes mirror
(capi.py) of the extern "C" library built from csrc/ that runs the step's four
phases as hand-written gfx950 HIP kernels.
"""
from . import capi  # noqa: F401
from . import configs  # noqa: F401
from . import deck  # noqa: F401
from . import driver  # noqa: F401
from . import grp  # noqa: F401
from . import layout  # noqa: F401
from . import learned  # noqa: F401
from . import mtb  # noqa: F401
from . import rsl  # noqa: F401
from . import timeloop  # noqa: F401
from . import tran  # noqa: F401
from . import unet  # noqa: F401
