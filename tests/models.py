"""Re-export of the package's model builders for the tests."""

from insem_amd.configs import *  # noqa: F401,F403
from insem_amd.configs import acoustic_box, grid, layered_coupled, quarter_annulus  # noqa: F401
