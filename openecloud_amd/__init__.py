"""openecloud_amd -- openECLOUD's per-timestep electron-cloud PIC loop on MI355X (gfx950).

Modules keep the reference's module and class names (grid.Grid, particles.Particles,
poissonSolver.PoissonSolver, particleBoundary.Absorb*, particleEmitter.FurmanEmitter /
homoLoader, particleManagement, magneticField, beam, randomGen); the compute runs in the HIP
kernels of liboecl.so.
"""
__all__ = ['grid', 'particles', 'poissonSolver', 'particleBoundary', 'particleEmitter',
           'particleManagement', 'magneticField', 'beam', 'randomGen', 'constants']
