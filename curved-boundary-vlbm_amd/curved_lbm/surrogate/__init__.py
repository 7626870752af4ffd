"""Learned warm start of the 3D solver: a small convolutional neural operator
(HIP BF16 MFMA kernels, include/vlbm_nn.h) trained on equilibria the solver
itself produces, whose prediction starts the solver near equilibrium."""
