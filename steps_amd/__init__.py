"""MI355X-native operator-split stochastic reaction-diffusion engine.

Drop-in for the STEPS TetOpSplit / DistTetOpSplit solvers: the model and mesh
objects mirror steps.model / steps.geom, ``steps_amd.solver`` holds the solver
classes, and the compute path is the HIP engine behind include/stg.h.
"""
__version__ = "0.1.0"
