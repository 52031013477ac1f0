"""openecloud_amd -- openECLOUD's per-timestep electron-cloud PIC loop on MI355X (gfx950).

Modules keep the reference's module and class names (grid.Grid, particles.Particles,
poissonSolver.PoissonSolver, particleBoundary.Absorb*, particleEmitter.FurmanEmitter /
homoLoader, particleManagement, magneticField, beam, randomGen); the compute runs in the HIP
kernels of liboecl.so.
"""
__all__ = ['grid', 'particles', 'poissonSolver', 'particleBoundary', 'particleEmitter',
           'particleManagement', 'magneticField', 'beam', 'randomGen', 'constants']

# set_host_arrays(True): the reference API's array getters return NumPy copies, as a reference
# driver script expects (numpy.sum / numpy.asarray / numpy.save on them).
from ._lib import host_arrays, set_host_arrays  # noqa: E402,F401
