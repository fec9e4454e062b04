"""MI355X island-model evolutionary-multitasking engine (HIP/gfx950)."""
from .engine import (DE_CURRENT_1_BIN, DE_RAND_1_BIN, DIAG, FUNC_IDS, GA_SBX_PM, ROTATE, SHIFT,  # noqa: F401
                     SCHEDULE_AUTO, SCHEDULE_EPOCHS, SCHEDULE_PERSISTENT,
                     FORMULAS_TEXTBOOK, FORMULAS_REFERENCE_DE, FORMULAS_REFERENCE_GA, ADMIT_AT_CREATE, ADMIT_ONLINE,
                     Engine)
from ._lib import EmtfError, LIB_PATH  # noqa: F401
from . import tasks  # noqa: F401
