"""Writes golden vectors of the reference's grid geometry.

Runs only in the build container, where the reference is mounted at /root/reference.  The
reference's grid.pyx and constants.pyx (GSL-free modules) are compiled with Cython into a temporary directory, imported, and evaluated on small grids;
inputs and outputs are saved to tests/golden/reference_grid_vectors.npz.  Nothing of the
reference is kept: the fixture holds arrays only.

These vectors pin what no closed form pins: the cut-cell bisection lengths, boundary points,
cut-cell point order and normals, and the edge-to-node rows (grid.pyx:468-1330).
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

REF = '/root/reference'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'reference_grid_vectors.npz')
# poissonSolver.pyx is not used: under the installed SciPy its eye(np, np, -nx) call
# receives -nx wrapped through an unsigned int and raises.  The Poisson solve is pinned by the
# linear system itself instead (tests/test_oracle_poisson.py).
MODULES = ['constants', 'grid']

SETUP = '''
from setuptools import setup, Extension
from Cython.Build import cythonize
import numpy
exts = [Extension(m, [m + '.pyx'], include_dirs=[numpy.get_include()],
                  extra_compile_args=['-O2', '-ffp-contract=off'])
        for m in %r]
setup(ext_modules=cythonize(exts, language_level=2, quiet=True))
''' % MODULES

CASES = [
    ('rect_16x12', (16, 12, 0.1, 0.05, 0)),
    ('ellip_20x17', (20, 17, 0.09, 0.07, 1)),
    ('ellip_33x29', (33, 29, 0.09, 0.07, 1)),
    ('circle_24x24', (24, 24, 0.04, 0.04, 1)),
]


def main():
    tmp = tempfile.mkdtemp(prefix='oecl_refgrid_')
    try:
        for m in MODULES:
            shutil.copy(os.path.join(REF, m + '.pyx'), tmp)
            if os.path.exists(os.path.join(REF, m + '.pxd')):
                shutil.copy(os.path.join(REF, m + '.pxd'), tmp)
        with open(os.path.join(tmp, 'setup_ref.py'), 'w') as f:
            f.write(SETUP)
        env = dict(os.environ, MPLBACKEND='Agg')
        subprocess.check_call([sys.executable, 'setup_ref.py', 'build_ext', '--inplace', '-q'],
                              cwd=tmp, env=env)
        sys.path.insert(0, tmp)
        os.environ['MPLBACKEND'] = 'Agg'
        import grid as rgrid
        out = {}
        for name, args in CASES:
            g = rgrid.Grid(*args)
            out[name + '/args'] = np.asarray(args, dtype=np.float64)
            for k in ('InsidePoints', 'InsideEdges', 'InsideFaces', 'Ds', 'Dsi', 'Da', 'Dai', 'Dst',
                      'Dat', 'BoundaryPoints', 'CutCellPointsInd', 'CutCellNormalVectors',
                      'CutCellCenter'):
                try:
                    out[name + '/' + k] = np.asarray(getattr(g, 'get' + k)())
                except Exception:
                    pass
            e2n = g.getEdgeToNode().tocoo()
            out[name + '/e2n_row'] = e2n.row.astype(np.int64)
            out[name + '/e2n_col'] = e2n.col.astype(np.int64)
            out[name + '/e2n_val'] = e2n.data.astype(np.float64)
        np.savez_compressed(OUT, **out)
        print('wrote', OUT, len(out), 'arrays')
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == '__main__':
    main()
