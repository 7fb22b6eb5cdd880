"""Configuration keys, defaults and numeric constants of the PythonMHD front-end
(Source/PythonMHD_Constants.py:7-79).  Init files written for the reference
use these names unchanged."""

SMALL_VALUE = 0.000000000001          # Source/PythonMHD_Constants.py:7
FOUR_THIRDS = 1.333333333333333       # Source/PythonMHD_Constants.py:8

OUTFLOW = 0
PERIODIC = 1

CARTESIAN = "Cartesian"
CYLINDRICAL = "Cylindrical"
SPHERICAL = "Spherical"

# gridPar keys
GRID_TYPE = "gridType"
NUM_X_CELLS = "numXCells"
MIN_X = "minX"
MAX_X = "maxX"
BC_X = "BcX"
NUM_Y_CELLS = "numYCells"
MIN_Y = "minY"
MAX_Y = "maxY"
BC_Y = "BcY"
NUM_Z_CELLS = "numZCells"
MIN_Z = "minZ"
MAX_Z = "maxZ"
BC_Z = "BcZ"
DEFAULT_BC_X = OUTFLOW
DEFAULT_BC_Y = OUTFLOW
DEFAULT_BC_Z = OUTFLOW

# simPar keys and defaults
IS_MHD = "isMHD"
TLIM = "tLim"
MAX_CYCLES = "maxCycles"
GAMMA = "gamma"
CFL = "cfl"
MIN_DENSITY = "minDens"
MIN_PRESSURE = "minPres"
MIN_ENERGY = "minEnergy"
RECONSTRUCT_ORDER = "reconstructOrder"
RIEMANN_SOLVER = "riemannSolver"
ENTROPY_FIX = "entropyFix"
EPSILON = "epsilon"
PLOT_DATA = "plotData"
SAVE_DATA = "saveData"

DEFAULT_IS_MHD = False
DEFAULT_MIN_DENSITY = 0.000000000001
DEFAULT_MIN_PRESSURE = 0.000000000001
DEFAULT_MIN_ENERGY = 0.000000000001
DEFAULT_CFL_1D = 0.8
DEFAULT_CFL_2D = 0.8
DEFAULT_CFL_3D = 0.4
DEFAULT_MAX_CYCLES = 50000
DEFAULT_GAMMA = 1.4
DEFAULT_EPSILON = 0.5

NO_SPATIAL_RECONSTRUCTION = 0
# PLM is commented out in the reference (Source/PythonMHD_Constants.py:77); this
# build adds it as order 1.
PLM_SPATIAL_RECONSTRUCTION = 1
PPM_SPATIAL_RECONSTRUCTION = 2
DEFAULT_SPATIAL_RECONSTRUCTION = PPM_SPATIAL_RECONSTRUCTION

# new in this build: arithmetic of the device step (64: FP64 throughout;
# 32: FP32 bulk with FP64 wave-speed/normalisation islands)
PRECISION = "precision"
DEFAULT_PRECISION = 64
