"""MI355X-native CTU+CT Godunov step for PythonMHD (drop-in for its evolve functions)."""
from . import constants  # noqa: F401
